// dmc_server.hip -- serving loop over the batched extraction (host code).
//
// Requests (one volume each) are queued by dmc_server_submit.  Each worker
// thread owns a DMC context (and so a stream of its own) and repeatedly
// takes the oldest waiting requests -- as many as fit the HBM budget and
// max_batch -- runs them as one dmc_extract_batch, copies the union mesh
// into a result block (one device-to-device copy per array) and completes
// the requests.  The passes of different workers overlap on the GPU: while
// one waits on a host round trip (a count read back between stages), the
// other's kernels run.  A request keeps its slice of the block until
// dmc_server_release; blocks are pooled and reused.  With max_batch = 1 and
// one worker the same loop serves one request at a time, the baseline the
// batched mode is measured against.
//
// Built only on the public C ABI of include/tmc_dmc.h plus HIP stream/copy
// calls; no counterpart in the reference (which has no serving path).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/tmc_dmc.h"

namespace {

using Clock = std::chrono::steady_clock;

// Upper bound of dmc_server_opts.workers (each worker holds a context's
// buffers, grown to its largest pass).
constexpr int kMaxWorkers = 8;
// Default bound of a pass's grid points (dmc_server_opts.max_batch_points
// = 0): keeps a pass's union well inside the engine's index limits (2^32
// active cells, 2^31 vertices, 2^29 quadrilaterals) for any mix of
// requests, each of which fits alone.
constexpr int64_t kDefaultBatchPoints = (int64_t)1 << 28;

// One pooled device allocation holding the meshes of one batch.
struct Block {
    void* p = nullptr;
    size_t cap = 0;
    int refs = 0;
};

struct Request {
    uint64_t ticket = 0;
    dmc_volume vol{};
    size_t est = 0;          // admission estimate (bytes)
    double t_submit = 0, t_start = 0, t_done = 0;
    int32_t batch_size = 0;
    dmc_status status = DMC_OK;
    bool done = false;
    dmc_mesh mesh{};         // device pointers into the block
    std::vector<int32_t> bnd;
    Block* block = nullptr;
    size_t held = 0;         // bytes of the block charged to this request
};

// Working set of one volume in the engine: the (row, segment) lists and
// counts (256 slots per row and 256-cell segment), plus a bound for the
// active-cell, mesh and halfedge arrays proportional to the points.
size_t estimate_bytes(const dmc_volume& v, int field_type, bool on_device) {
    const size_t nx = (size_t)v.grid.nx, ny = (size_t)v.grid.ny, nz = (size_t)v.grid.nz;
    const size_t rows = (ny - 1) * (nz - 1), segs = (nx - 1 + 255) / 256;
    const size_t pts = nx * ny * nz;
    size_t b = rows * segs * (256 * 4 + 8) + 40 * pts;
    if (!on_device) b += pts * (field_type == DMC_FIELD_F32 ? 4 : 8);
    return b;
}

}  // namespace

// One worker: a context and a stream for its result copies.
struct Worker {
    dmc_context* ctx = nullptr;
    hipStream_t copy_stream = nullptr;
    std::thread th;
    size_t pool_bytes = 0;   // its context's buffer pools after its last pass
};

struct dmc_server {
    dmc_server_opts opts{};
    std::vector<std::unique_ptr<Worker>> workers;
    hipStream_t dl_stream = nullptr;    // downloads (caller threads, one at a time)
    std::mutex mu, dl_mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request*> queue;
    std::map<uint64_t, std::unique_ptr<Request>> reqs;
    std::vector<Block*> blocks;         // every block (pooled)
    size_t held_bytes = 0;              // result blocks holding unreleased results (their full capacity)
    size_t idle_bytes = 0;              // result blocks free for reuse
    size_t ctx_bytes = 0;               // buffer pools the worker contexts keep between passes
    size_t inflight_bytes = 0;          // estimated working sets of the passes running
    uint64_t next_ticket = 1;
    bool stop = false;
    Clock::time_point t0 = Clock::now();
    std::string err;
    // totals
    uint64_t batches = 0, served = 0;

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
};

namespace {

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Bytes of a block for a union mesh of V vertices and Q quadrilaterals:
// positions, normals, quads, triangles, twins, halfedge flags, he_v.
size_t block_bytes(int64_t V, int64_t Q) {
    return align256(24 * V) * 2 + align256(16 * Q) + align256(24 * Q) + align256(16 * Q) + align256(4 * Q) +
           align256(4 * V);
}

// A free block of at least `bytes` (smallest fit), or a new one.  Called
// with the lock held; the block leaves the idle pool.
Block* get_block(dmc_server* s, size_t bytes) {
    Block* best = nullptr;
    for (Block* b : s->blocks)
        if (b->refs == 0 && b->cap >= bytes && (!best || b->cap < best->cap)) best = b;
    if (best) {
        s->idle_bytes -= best->cap;
        return best;
    }
    Block* b = new Block();
    if (hipMalloc(&b->p, bytes) != hipSuccess) {
        (void)hipGetLastError();
        delete b;
        return nullptr;
    }
    b->cap = bytes;
    s->blocks.push_back(b);
    return b;
}

// Free the idle result blocks (lock held).
void free_idle_blocks(dmc_server* s) {
    std::vector<Block*> keep;
    for (Block* b : s->blocks) {
        if (b->refs == 0) {
            (void)hipFree(b->p);
            delete b;
        } else {
            keep.push_back(b);
        }
    }
    s->blocks.swap(keep);
    s->idle_bytes = 0;
}

// Device memory the server holds or is about to use (lock held): result
// blocks (held and idle), the worker contexts' pools and the estimated
// working sets of the passes running.
size_t committed(const dmc_server* s) { return s->held_bytes + s->idle_bytes + s->ctx_bytes + s->inflight_bytes; }

// Whether `est` more bytes fit the budget, freeing idle blocks if that
// makes them fit (lock held).
bool fits(dmc_server* s, size_t est) {
    const size_t budget = s->opts.budget_bytes;
    if (est + committed(s) <= budget) return true;
    if (s->idle_bytes > 0 && est + committed(s) - s->idle_bytes <= budget) {
        free_idle_blocks(s);
        return true;
    }
    return false;
}

// One pass over `batch`: its status (every member gets it).  With
// complete = false a failed pass leaves its members open (the caller
// retries them one by one).
dmc_status run_pass(dmc_server* s, Worker* w, std::vector<Request*>& batch, bool complete) {
    const int n = (int)batch.size();
    std::vector<dmc_volume> vols(n);
    for (int i = 0; i < n; i++) vols[i] = batch[i]->vol;
    std::vector<dmc_mesh> out(n);
    // pools that would keep the server over its budget (a context grown by
    // an earlier, larger pass) are freed first: the pass then allocates only
    // what it needs
    bool trim = false;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        trim = w->pool_bytes > 0 && committed(s) > s->opts.budget_bytes;
    }
    if (trim && dmc_context_trim(w->ctx) == DMC_OK) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->ctx_bytes -= w->pool_bytes;
        w->pool_bytes = 0;
    }
    const double t_start = s->now();
    dmc_status st = dmc_extract_batch(w->ctx, vols.data(), n, s->opts.field_type, s->opts.flags, out.data());
    int64_t V = 0, Q = 0;
    for (int i = 0; i < n && st == DMC_OK; i++) {
        V += out[i].nv;
        Q += out[i].nq;
    }
    Block* blk = nullptr;
    std::vector<std::vector<int32_t>> bnd(n);
    if (st == DMC_OK) {
        for (int i = 0; i < n; i++) {
            bnd[i].resize(out[i].nbnd);
            if (out[i].nbnd)
                st = dmc_batch_download(w->ctx, i, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                                        bnd[i].data());
            if (st != DMC_OK) break;
        }
    }
    if (st == DMC_OK) {
        std::lock_guard<std::mutex> lk(s->mu);
        blk = get_block(s, std::max<size_t>(block_bytes(V, Q), 256));
        if (!blk) st = DMC_ERR_OUT_OF_MEMORY;
        else blk->refs = n;
    }
    if (st == DMC_OK && (V > 0 || Q > 0)) {
        // the union arrays are contiguous in volume order, starting at
        // volume 0's pointers
        char* p = static_cast<char*>(blk->p);
        const dmc_mesh& m0 = out[0];
        const bool tris = (s->opts.flags & DMC_TRIANGLES) != 0;
        struct Sec {
            const void* src;
            size_t bytes;
            size_t off;
        };
        size_t off = 0;
        std::vector<Sec> secs;
        auto add = [&](const void* src, size_t bytes, size_t room) {
            secs.push_back(Sec{src, bytes, off});
            off += align256(room);
        };
        add(m0.d_positions, 24 * V, 24 * V);
        add(m0.d_normals, 24 * V, 24 * V);
        add(m0.d_quads, 16 * Q, 16 * Q);
        add(tris ? m0.d_tris : nullptr, tris ? 24 * Q : 0, 24 * Q);
        add(m0.d_twins, 16 * Q, 16 * Q);
        add(m0.d_he_flags, 4 * Q, 4 * Q);
        add(m0.d_he_v, 4 * V, 4 * V);
        for (const Sec& x : secs)
            if (x.src && x.bytes && st == DMC_OK &&
                hipMemcpyAsync(p + x.off, x.src, x.bytes, hipMemcpyDeviceToDevice, w->copy_stream) != hipSuccess)
                st = DMC_ERR_HIP;
        if (hipStreamSynchronize(w->copy_stream) != hipSuccess) st = DMC_ERR_HIP;
        // re-point every volume's mesh into the block
        int64_t vb = 0, qb = 0;
        for (int i = 0; i < n; i++) {
            dmc_mesh& m = out[i];
            m.d_positions = reinterpret_cast<const double*>(p + secs[0].off) + 3 * vb;
            m.d_normals = reinterpret_cast<const double*>(p + secs[1].off) + 3 * vb;
            m.d_quads = reinterpret_cast<const int32_t*>(p + secs[2].off) + 4 * qb;
            m.d_tris = tris ? reinterpret_cast<const int32_t*>(p + secs[3].off) + 6 * qb : nullptr;
            m.d_twins = reinterpret_cast<const int32_t*>(p + secs[4].off) + 4 * qb;
            m.d_he_flags = reinterpret_cast<const uint8_t*>(p + secs[5].off) + 4 * qb;
            m.d_he_v = reinterpret_cast<const int32_t*>(p + secs[6].off) + vb;
            vb += m.nv;
            qb += m.nq;
        }
    }
    const double t_done = s->now();
    size_t cb = 0;
    (void)dmc_context_bytes(w->ctx, &cb);
    std::lock_guard<std::mutex> lk(s->mu);
    s->batches++;
    s->ctx_bytes += cb;
    s->ctx_bytes -= w->pool_bytes;
    w->pool_bytes = cb;
    if (st != DMC_OK) {
        s->err = dmc_last_error(w->ctx);
        if (blk) {
            blk->refs = 0;
            s->idle_bytes += blk->cap;
        }
        if (!complete) return st;
    }
    // each member is charged its share of the block's whole capacity
    size_t left = blk ? blk->cap : 0;
    const int64_t tot = std::max<int64_t>(1, V + Q);
    for (int i = 0; i < n; i++) {
        Request* r = batch[i];
        r->t_start = t_start;
        r->t_done = t_done;
        r->batch_size = n;
        r->status = st;
        if (st == DMC_OK) {
            r->mesh = out[i];
            r->bnd = std::move(bnd[i]);
            r->block = blk;
            r->held = i + 1 == n ? left : (size_t)((double)blk->cap * (double)(out[i].nv + out[i].nq) / (double)tot);
            r->held = std::min(r->held, left);
            left -= r->held;
            s->held_bytes += r->held;
        } else {
            r->mesh = dmc_mesh{};
        }
        r->done = true;
        s->served++;
    }
    s->cv_done.notify_all();
    return st;
}

// A pass over the admitted requests.  A batch that fails on a size limit of
// the union (DMC_ERR_CAPACITY) or on one bad member (DMC_ERR_INVALID_ARG)
// is run again one request at a time, so that only the requests that fail
// alone fail.
void run_batch(dmc_server* s, Worker* w, std::vector<Request*>& batch) {
    const bool retry = batch.size() > 1;
    const dmc_status st = run_pass(s, w, batch, !retry);
    if (st == DMC_OK || !retry) return;
    if (st != DMC_ERR_CAPACITY && st != DMC_ERR_INVALID_ARG) {
        // complete every member with the pass's status
        std::vector<Request*> none;
        std::lock_guard<std::mutex> lk(s->mu);
        for (Request* r : batch) {
            r->t_start = r->t_done = s->now();
            r->batch_size = (int32_t)batch.size();
            r->status = st;
            r->mesh = dmc_mesh{};
            r->done = true;
            s->served++;
        }
        s->cv_done.notify_all();
        return;
    }
    for (Request* r : batch) {
        std::vector<Request*> one{r};
        run_pass(s, w, one, true);
    }
}

void worker_main(dmc_server* s, Worker* w) {
    (void)hipSetDevice(s->opts.device);
    const size_t budget = s->opts.budget_bytes;
    // a pass may start when no other is running, or when the oldest request
    // fits beside what the server already holds (one that can never fit is
    // failed at once)
    auto ready = [&] {
        if (s->stop) return true;
        if (s->queue.empty()) return false;
        const size_t est = s->queue.front()->est;
        return s->inflight_bytes == 0 || est > budget || fits(s, est);
    };
    for (;;) {
        std::vector<Request*> batch;
        size_t used = 0;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_work.wait(lk, ready);
            if (s->stop && s->queue.empty()) return;
            // admission: oldest first, while the working set fits beside what
            // the server holds (results, idle blocks, worker pools, passes in
            // flight)
            const int64_t max_points = s->opts.max_batch_points > 0 ? s->opts.max_batch_points : kDefaultBatchPoints;
            int64_t points = 0;
            while (!s->queue.empty() && (int)batch.size() < s->opts.max_batch) {
                Request* r = s->queue.front();
                if (r->est > budget) {
                    s->queue.pop_front();
                    r->status = DMC_ERR_BUDGET;
                    r->done = true;
                    r->t_start = r->t_done = s->now();
                    s->cv_done.notify_all();
                    continue;
                }
                const int64_t pts = (int64_t)r->vol.grid.nx * r->vol.grid.ny * r->vol.grid.nz;
                if (!batch.empty() && !fits(s, used + r->est)) break;
                if (!batch.empty() && points + pts > max_points) break;
                batch.push_back(r);
                used += r->est;
                points += pts;
                s->queue.pop_front();
            }
            s->inflight_bytes += used;
            // requests left over are for another idle worker
            if (!s->queue.empty()) s->cv_work.notify_one();
        }
        if (!batch.empty()) run_batch(s, w, batch);
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->inflight_bytes -= used;
        }
        s->cv_work.notify_all();
    }
}

}  // namespace

extern "C" {

dmc_status dmc_server_create(dmc_server** out, const dmc_server_opts* opts) {
    if (!out || !opts) return DMC_ERR_INVALID_ARG;
    *out = nullptr;
    if (opts->max_batch < 1 || opts->workers < 0 || opts->workers > kMaxWorkers ||
        (opts->field_type != DMC_FIELD_F32 && opts->field_type != DMC_FIELD_F64))
        return DMC_ERR_INVALID_ARG;
    dmc_server* s = new dmc_server();
    s->opts = *opts;
    if (s->opts.budget_bytes == 0) s->opts.budget_bytes = (size_t)96 << 30;
    if (s->opts.workers == 0) s->opts.workers = 1;
    s->opts.flags &= ~(uint32_t)DMC_STAGE_TIMING;
    dmc_status st = DMC_OK;
    if (hipSetDevice(opts->device) != hipSuccess ||
        hipStreamCreateWithFlags(&s->dl_stream, hipStreamNonBlocking) != hipSuccess)
        st = DMC_ERR_HIP;
    for (int i = 0; i < s->opts.workers && st == DMC_OK; i++) {
        auto w = std::make_unique<Worker>();
        st = dmc_create(&w->ctx, opts->device);
        if (st == DMC_OK && hipStreamCreateWithFlags(&w->copy_stream, hipStreamNonBlocking) != hipSuccess)
            st = DMC_ERR_HIP;
        s->workers.push_back(std::move(w));
    }
    if (st != DMC_OK) {
        (void)hipGetLastError();
        for (auto& w : s->workers) {
            if (w->copy_stream) (void)hipStreamDestroy(w->copy_stream);
            if (w->ctx) dmc_destroy(w->ctx);
        }
        if (s->dl_stream) (void)hipStreamDestroy(s->dl_stream);
        delete s;
        return st;
    }
    for (auto& w : s->workers) w->th = std::thread(worker_main, s, w.get());
    *out = s;
    return DMC_OK;
}

dmc_status dmc_server_destroy(dmc_server* s) {
    if (!s) return DMC_ERR_INVALID_ARG;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto& w : s->workers)
        if (w->th.joinable()) w->th.join();
    (void)hipSetDevice(s->opts.device);
    for (auto& w : s->workers) (void)hipStreamSynchronize(w->copy_stream);
    for (Block* b : s->blocks) {
        if (b->p) (void)hipFree(b->p);
        delete b;
    }
    (void)hipStreamSynchronize(s->dl_stream);
    for (auto& w : s->workers) {
        (void)hipStreamDestroy(w->copy_stream);
        dmc_destroy(w->ctx);
    }
    (void)hipStreamDestroy(s->dl_stream);
    delete s;
    return DMC_OK;
}

dmc_status dmc_server_submit(dmc_server* s, const dmc_volume* vol, uint64_t* ticket) {
    if (!s || !vol || !ticket || !vol->field) return DMC_ERR_INVALID_ARG;
    // the engine's grid rules, checked here so that a request that can
    // never run fails alone instead of failing the pass it would join
    const dmc_grid& g = vol->grid;
    const char* bad = nullptr;
    dmc_status st = DMC_ERR_INVALID_ARG;
    if (g.nx < 2 || g.ny < 2 || g.nz < 2) bad = "grid needs at least 2 points per axis";
    else if (g.nx > (1 << 24) || g.ny > (1 << 24) || g.nz > (1 << 24)) bad = "an axis above 16777216 points";
    else if ((uint64_t)(g.nx - 1) * (uint64_t)(g.ny - 1) * (uint64_t)(g.nz - 1) > 0xFFFFFFFFull) {
        bad = "more than 2^32 cells";
        st = DMC_ERR_CAPACITY;
    } else if ((s->opts.flags & DMC_GRADIENTS_SUPPLIED) && !vol->gradients) {
        bad = "gradients are NULL";
    }
    if (bad) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->err = bad;
        return st;
    }
    auto r = std::make_unique<Request>();
    r->vol = *vol;
    r->est = estimate_bytes(*vol, s->opts.field_type, (s->opts.flags & DMC_FIELD_ON_DEVICE) != 0);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        r->ticket = s->next_ticket++;
        r->t_submit = s->now();
        *ticket = r->ticket;
        s->queue.push_back(r.get());
        s->reqs.emplace(r->ticket, std::move(r));
    }
    s->cv_work.notify_one();
    return DMC_OK;
}

dmc_status dmc_server_poll(dmc_server* s, uint64_t ticket, int* done) {
    if (!s || !done) return DMC_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return DMC_ERR_INVALID_ARG;
    *done = it->second->done ? 1 : 0;
    return DMC_OK;
}

dmc_status dmc_server_wait(dmc_server* s, uint64_t ticket, dmc_result* out) {
    if (!s) return DMC_ERR_INVALID_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return DMC_ERR_INVALID_ARG;
    Request* r = it->second.get();
    s->cv_done.wait(lk, [&] { return r->done; });
    if (out) {
        out->mesh = r->mesh;
        out->t_submit = r->t_submit;
        out->t_start = r->t_start;
        out->t_done = r->t_done;
        out->batch_size = r->batch_size;
        out->status = r->status;
    }
    return r->status;
}

dmc_status dmc_server_download(dmc_server* s, uint64_t ticket, double* positions, double* normals, int32_t* quads,
                               int32_t* tris, int32_t* halfedges, int32_t* he_v, int32_t* he_f, int32_t* bnd) {
    if (!s) return DMC_ERR_INVALID_ARG;
    dmc_mesh m;
    std::vector<int32_t> b;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->reqs.find(ticket);
        if (it == s->reqs.end()) return DMC_ERR_INVALID_ARG;
        Request* r = it->second.get();
        if (!r->done) return DMC_ERR_STATE;
        if (r->status != DMC_OK) return r->status;
        m = r->mesh;
        b = r->bnd;
    }
    (void)hipSetDevice(s->opts.device);
    // the request's slice of its block stays valid until it is released
    std::lock_guard<std::mutex> dl(s->dl_mu);
    auto d2h = [&](void* dst, const void* src, size_t bytes) {
        return (dst && bytes) ? hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s->dl_stream) : hipSuccess;
    };
    hipError_t e = hipSuccess;
    std::vector<int32_t> q, tw;
    std::vector<uint8_t> fl;
    if (halfedges) {
        q.resize(4 * m.nq);
        tw.resize(4 * m.nq);
        fl.resize(4 * m.nq);
    }
    const hipError_t es[] = {d2h(positions, m.d_positions, 24 * m.nv), d2h(normals, m.d_normals, 24 * m.nv),
                             d2h(quads, m.d_quads, 16 * m.nq), d2h(tris, m.d_tris, 12 * m.nt),
                             d2h(he_v, m.d_he_v, 4 * m.nv),
                             d2h(halfedges ? q.data() : nullptr, m.d_quads, 16 * m.nq),
                             d2h(halfedges ? tw.data() : nullptr, m.d_twins, 16 * m.nq),
                             d2h(halfedges ? fl.data() : nullptr, m.d_he_flags, 4 * m.nq),
                             hipStreamSynchronize(s->dl_stream)};
    for (hipError_t x : es)
        if (x != hipSuccess) e = x;
    if (e != hipSuccess) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->err = hipGetErrorString(e);
        return DMC_ERR_HIP;
    }
    if (halfedges)
        for (int64_t h = 0; h < 4 * m.nq; h++) {
            int32_t* o = halfedges + 5 * h;
            o[0] = q[h];
            o[1] = (int32_t)(h / 4);
            o[2] = (int32_t)((h & ~3) | ((h + 1) & 3));
            o[3] = tw[h];
            o[4] = fl[h];
        }
    if (he_f)
        for (int64_t f = 0; f < m.nq; f++) he_f[f] = (int32_t)(4 * f);
    if (bnd && !b.empty()) memcpy(bnd, b.data(), b.size() * sizeof(int32_t));
    return DMC_OK;
}

dmc_status dmc_server_release(dmc_server* s, uint64_t ticket) {
    if (!s) return DMC_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return DMC_ERR_INVALID_ARG;
    Request* r = it->second.get();
    if (!r->done) return DMC_ERR_STATE;
    if (r->block && --r->block->refs == 0) s->idle_bytes += r->block->cap;
    s->held_bytes -= r->held;
    s->reqs.erase(it);
    return DMC_OK;
}

dmc_status dmc_server_stats(dmc_server* s, uint64_t* batches, uint64_t* served, size_t* held_bytes,
                            size_t* pool_bytes) {
    if (!s) return DMC_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    size_t pool = s->ctx_bytes;
    for (Block* b : s->blocks) pool += b->cap;
    if (batches) *batches = s->batches;
    if (served) *served = s->served;
    if (held_bytes) *held_bytes = s->held_bytes;
    if (pool_bytes) *pool_bytes = pool;
    return DMC_OK;
}

const char* dmc_server_last_error(dmc_server* s) { return s ? s->err.c_str() : "null server"; }

}  // extern "C"

// bemx_serve.hip -- device-resident request loop for BEM solves on one GPU.
//
// Continuous batching: requests (a mesh, a material, a frequency and 1..4
// load cases) are queued as they arrive.  NSLOTS worker threads each own a
// device context (its own streams and grow-only workspace); a free worker
// takes the oldest queued request as soon as the HBM it needs fits in what
// the budget has left, together with up to batch - 1 more queued requests
// that can share its launches (the same mesh and KODE and as many load
// cases), runs assembly + factorisation + refinement for them as one batch
// (ctx_solve_items; every launch of the mixed factorisation and refinement
// covers the whole batch) and hands the solutions back.  Requests that arrive while
// others are on the GPU are admitted as slots and memory free up, and the
// work of different slots overlaps on the device: while one solve sits in
// the latency-bound chain of its factorisation, another's bulk update fills
// the CUs.  A slot keeps the mesh and static pass (GHMATECE) of the last
// mesh it served, so a stream of requests on one mesh (a frequency sweep)
// pays the static pass once per slot, and it keeps one captured
// factorisation graph per size class.  Each request's arithmetic is the
// same whichever slot runs it and whatever runs beside it.
//
// Admission charges each slot what its workspace can hold: the bound of
// ctx_solve_bytes_bound at the largest sizes the slot has served (its
// buffers only grow), not the footprint of the request it runs.  The sum of
// the charges never exceeds the budget.  When the oldest request does not
// fit, idle slots give their workspace back (the charge drops to zero) until
// it does.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <condition_variable>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/bemx.h"
#include "bemx_lu.hpp"

namespace bemx {
size_t ctx_solve_bytes_bound(int n, int nbe, int np, int npg, int nrhs, int mode, int nb);
int ctx_solve_items(bemx_ctx* c, int nitem, SolveItem* items, int mode,
                    const std::function<void(int, int)>* batch_done);
int batch_size();
size_t ctx_device_bytes(bemx_ctx* c);
void ctx_release_workspace(bemx_ctx* c);
void ctx_graph_counts(const bemx_ctx* c, long long* captures, long long* replays,
                      long long* precaptures);
void ctx_capture_ahead(bemx_ctx* c, int ahead);
}  // namespace bemx

namespace {

using clk = std::chrono::steady_clock;

double ms_between(clk::time_point a, clk::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

uint64_t fnv1a(uint64_t h, const void* p, size_t n) {
    const unsigned char* b = static_cast<const unsigned char*>(p);
    for (size_t i = 0; i < n; ++i) {
        h ^= b[i];
        h *= 1099511628211ull;
    }
    return h;
}

struct Request {
    long long id = 0;
    std::vector<double> cx, cy, cz;
    std::vector<int> cone;
    int np = 0, n = 0, nbe = 0, npg = 0;
    double ge = 0, rnu = 0, dam = 0, rho = 0, fr = 0;
    int nrhs = 1;
    std::vector<double> zdfi;  // interleaved complex, 3NBE x nrhs
    std::vector<int> kode;     // 3NBE x kode_cols
    int kode_cols = 1;         // 1: one KODE for every case; nrhs: one per case
    uint64_t mesh_key = 0;
    size_t footprint = 0;
    clk::time_point t_submit, t_start, t_done;
    bool done = false;
    int status = BEMX_OK;
    std::string err;
    std::vector<double> x;
    double stats[4] = {0, 0, 0, 0};
    int slot = -1;
    int batch = 1;  // requests in the batch it was solved in
};

// Largest sizes a slot has served: its buffers hold (at most) the bound at
// these sizes.
struct SlotSizes {
    int n = 0, nbe = 0, np = 0, npg = 0, nrhs = 0, nb = 0;
    SlotSizes grown(const Request& r, int batch) const {
        SlotSizes g;
        g.n = std::max(n, r.n);
        g.nbe = std::max(nbe, r.nbe);
        g.np = std::max(np, r.np);
        g.npg = std::max(npg, r.npg);
        g.nrhs = std::max(nrhs, r.nrhs);
        g.nb = std::max(nb, batch);
        return g;
    }
};

size_t bound_of(const SlotSizes& z, int mode) {
    if (!z.nbe) return 0;
    return bemx::ctx_solve_bytes_bound(z.n, z.nbe, z.np, z.npg, z.nrhs, mode, z.nb);
}

// Whether request b can share the launches of a batch led by request a: the
// same mesh and static pass, one KODE for every load case, the same KODE
// and the same number of load cases (the load cases themselves may differ).
bool batchable(const Request& a, const Request& b) {
    return a.kode_cols == 1 && b.kode_cols == 1 && a.mesh_key == b.mesh_key &&
           a.nrhs == b.nrhs && a.kode == b.kode;
}

struct Slot {
    bemx_ctx* ctx = nullptr;
    SlotSizes sizes;
    size_t charge = 0;  // bytes charged against the budget
    bool busy = false;
};

}  // namespace

struct bemx_server {
    int device = 0;
    clk::time_point t0 = clk::now();
    int mode = BEMX_MODE_MIXED;
    size_t budget = 0, charged = 0, peak_charged = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request*> queue;
    std::map<long long, std::unique_ptr<Request>> reqs;
    long long next_id = 1;
    int batch = 1;  // most requests per batch
    bool stop = false;
    std::vector<std::thread> workers;
    std::vector<Slot> slots;
    // statistics
    long long completed = 0, releases = 0, over_charge = 0, batches = 0;
    long long captures = 0, replays = 0, precaptures = 0;
    std::vector<long long> slot_captures, slot_replays, slot_precaptures;
    std::string err;
};

namespace {

size_t request_footprint(const Request& r, int mode) {
    return bemx::ctx_solve_bytes_bound(r.n, r.nbe, r.np, r.npg, r.nrhs, mode, 1);
}

// The mesh and static pass of request r on the slot's context (kept from
// the last request when it is the same mesh).
int ensure_mesh(bemx_ctx* c, uint64_t* slot_key, const Request* r) {
    int st = BEMX_OK;
    if (*slot_key != r->mesh_key) {
        *slot_key = 0;
        st = bemx_send_shared_data(c, r->cx.data(), r->cy.data(), r->cz.data(), nullptr, nullptr,
                                   nullptr, nullptr, nullptr, nullptr, r->cone.data(), r->np,
                                   r->npg, r->n, r->nbe);
        if (!st)
            st = bemx_ghmatece(c, r->ge, r->rnu, nullptr, nullptr, nullptr, nullptr, nullptr,
                               nullptr, nullptr, nullptr);
        if (!st) *slot_key = r->mesh_key;
    }
    return st;
}

void run_request(bemx_server* s, bemx_ctx* c, uint64_t* slot_key, Request* r) {
    int st = ensure_mesh(c, slot_key, r);
    double zge[2], zcs[2], zcp[2];
    if (!st) st = bemx_material(r->ge, r->rnu, r->dam, r->rho, zge, zcs, zcp);
    if (!st)
        st = r->kode_cols == 1
                 ? bemx_send_load_cases(c, r->nrhs, r->zdfi.data(), r->kode.data())
                 : bemx_send_load_cases_kode(c, r->nrhs, r->zdfi.data(), r->kode.data());
    if (!st) {
        r->x.assign((size_t)6 * r->nbe * r->nrhs, 0.0);
        st = bemx_solve(c, zge, zcs, zcp, r->fr, s->mode, r->x.data(), r->stats);
    }
    r->status = st;
    if (st) r->err = bemx_last_error(c);
}

void finish(bemx_server* s, Request* r);
void count_graphs(bemx_server* s, bemx_ctx* c, int slot);

// Requests batchable with the first: one upload of the KODE, every
// request's load cases, and the batched solves (up to two batches, the
// second one's assembly beside the first one's solve); each batch's
// requests are handed back as soon as it is done.
void run_batch(bemx_server* s, int slot, bemx_ctx* c, uint64_t* slot_key,
               std::vector<Request*>& rs) {
    Request* r0 = rs[0];
    int st = ensure_mesh(c, slot_key, r0);
    if (!st) st = bemx_send_load_cases(c, r0->nrhs, r0->zdfi.data(), r0->kode.data());
    std::vector<bemx::SolveItem> items(rs.size());
    for (size_t i = 0; i < rs.size() && !st; ++i) {
        Request* r = rs[i];
        bemx::SolveItem& it = items[i];
        st = bemx_material(r->ge, r->rnu, r->dam, r->rho, it.zge, it.zcs, it.zcp);
        it.fr = r->fr;
        it.zdfi = r->zdfi.data();
        r->x.assign((size_t)6 * r->nbe * r->nrhs, 0.0);
        it.x = r->x.data();
    }
    // a request handed back belongs to its waiter, who may free it: it is
    // not touched again (finished[] is this thread's record)
    std::vector<char> finished(rs.size(), 0);
    const std::function<void(int, int)> done = [&](int f0, int nf) {
        std::lock_guard<std::mutex> lk(s->mu);
        ++s->batches;
        count_graphs(s, c, slot);  // before the batch is handed back
        const clk::time_point now = clk::now();
        for (int i = f0; i < f0 + nf; ++i) {
            for (int q = 0; q < 4; ++q) rs[i]->stats[q] = items[i].stats[q];
            rs[i]->status = BEMX_OK;
            rs[i]->batch = nf;
            finished[i] = 1;
            finish(s, rs[i]);
        }
        // the next batch's solve starts now: its service time runs from here
        for (size_t i = f0 + nf; i < rs.size() && i < (size_t)(f0 + 2 * nf); ++i)
            rs[i]->t_start = now;
    };
    if (!st) st = bemx::ctx_solve_items(c, (int)rs.size(), items.data(), s->mode, &done);
    std::lock_guard<std::mutex> lk(s->mu);
    for (size_t i = 0; i < rs.size(); ++i) {
        if (finished[i]) continue;
        rs[i]->status = st ? st : BEMX_ERR_STATE;
        rs[i]->err = st ? bemx_last_error(c) : "request left unsolved by its batch";
        finish(s, rs[i]);
    }
}

// Adds the graph captures, replays and captures ahead of the slot's context
// since the last count to the server's statistics (the server's lock held).
void count_graphs(bemx_server* s, bemx_ctx* c, int slot) {
    long long cap = 0, rep = 0, pre = 0;
    bemx::ctx_graph_counts(c, &cap, &rep, &pre);
    s->captures += cap - s->slot_captures[slot];
    s->replays += rep - s->slot_replays[slot];
    s->precaptures += pre - s->slot_precaptures[slot];
    s->slot_captures[slot] = cap;
    s->slot_replays[slot] = rep;
    s->slot_precaptures[slot] = pre;
}

// Hands request r back to its waiter (the server's lock held).
void finish(bemx_server* s, Request* r) {
    r->t_done = clk::now();
    r->done = true;
    ++s->completed;
    s->cv_done.notify_all();
}

// Most batches one worker takes at once of requests batchable with the
// oldest one (BEMX_TAKE_BATCHES, 1..64; default 16).
int take_batches() {
    static const int t = [] {
        const char* v = getenv("BEMX_TAKE_BATCHES");
        const int x = v ? atoi(v) : 16;
        return std::min(64, std::max(1, x));
    }();
    return t;
}

void worker_main(bemx_server* s, int slot) {
    bemx_ctx* c = s->slots[slot].ctx;
    uint64_t key = 0;
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        Slot& me = s->slots[slot];
        // FIFO admission: the oldest request runs as soon as a slot is free
        // and the slot's charge, grown to that request (and to a batch when
        // the request can lead one and the budget has room for it), fits in
        // the budget
        bool give_back = false;
        int nb = 1;  // device slots for the requests taken (two batches)
        s->cv_work.wait(lk, [&] {
            if (s->stop) return true;
            if (s->queue.empty()) return false;
            const Request& f = *s->queue.front();
            nb = (s->batch > 1 && f.kode_cols == 1) ? 2 * s->batch : 1;
            for (;; nb = 1) {
                const size_t need = bound_of(me.sizes.grown(f, nb), s->mode);
                if (s->charged - me.charge + need <= s->budget) return true;
                if (nb == 1) break;
            }
            give_back = me.charge > 0;  // the request waits for memory: return ours
            return give_back;
        });
        if (s->stop) break;
        if (give_back) {
            lk.unlock();
            bemx::ctx_release_workspace(c);
            key = 0;
            lk.lock();
            s->charged -= me.charge;
            me.charge = 0;
            me.sizes = SlotSizes();
            ++s->releases;
            s->cv_work.notify_all();
            continue;
        }
        std::vector<Request*> rs(1, s->queue.front());
        s->queue.pop_front();
        // the queued requests that can share its launches, oldest first: up
        // to take_batches() batches, solved in turn through two batches of
        // device slots (each batch's assembly beside the previous batch's
        // solve) and handed back batch by batch
        const int take = nb > 1 ? take_batches() * s->batch : 1;
        for (auto it = s->queue.begin(); it != s->queue.end() && (int)rs.size() < take;) {
            if (batchable(*rs[0], **it)) {
                rs.push_back(*it);
                it = s->queue.erase(it);
            } else {
                ++it;
            }
        }
        me.sizes = me.sizes.grown(*rs[0], std::min(nb, (int)rs.size()));
        const size_t need = bound_of(me.sizes, s->mode);
        s->charged = s->charged - me.charge + need;
        me.charge = need;
        s->peak_charged = std::max(s->peak_charged, s->charged);
        me.busy = true;
        const clk::time_point t_start = clk::now();
        for (Request* r : rs) {
            r->slot = slot;
            r->t_start = t_start;
        }
        if (rs.size() == 1) ++s->batches;
        lk.unlock();
        if (rs.size() == 1)
            run_request(s, c, &key, rs[0]);
        else
            run_batch(s, slot, c, &key, rs);
        const size_t held = bemx::ctx_device_bytes(c);
        lk.lock();
        if (held > me.charge) ++s->over_charge;  // the bound was too small
        count_graphs(s, c, slot);
        me.busy = false;
        if (rs.size() == 1) finish(s, rs[0]);  // run_batch hands back its own
        s->cv_work.notify_all();
    }
    lk.unlock();
    // The objects this thread created for the context's solves (its second
    // stream, events and captured graphs) and the workspace are released by
    // this thread, after its last work has finished.
    bemx::ctx_release_workspace(c);
}

}  // namespace

extern "C" {

int bemx_material(double ge, double rnu, double dam, double rho, double* zge, double* zcs,
                  double* zcp) {
    // mu* = G (1 + 2i DAM), cs = sqrt(mu*/rho), cp = cs sqrt((2-2nu)/(1-2nu))
    if (!zge || !zcs || !zcp || !(rho > 0.0) || rnu >= 0.5) return BEMX_ERR_ARG;
    const std::complex<double> g(ge, ge * 2.0 * dam);
    const std::complex<double> cs = std::sqrt(g / rho);
    const std::complex<double> cp = cs * std::sqrt((2.0 - 2.0 * rnu) / (1.0 - 2.0 * rnu));
    zge[0] = g.real();
    zge[1] = g.imag();
    zcs[0] = cs.real();
    zcs[1] = cs.imag();
    zcp[0] = cp.real();
    zcp[1] = cp.imag();
    return BEMX_OK;
}

double bemx_request_footprint(int nbe, int n, int np, int npg, int nrhs, int mode) {
    if (nbe <= 0 || n < nbe || np <= 0 || npg <= 0 || nrhs < 1) return 0.0;
    if (mode != BEMX_MODE_FP64 && mode != BEMX_MODE_MIXED) return 0.0;
    return (double)bemx::ctx_solve_bytes_bound(n, nbe, np, npg, nrhs, mode, 1);
}

int bemx_server_create(int device, int nslots, double hbm_budget_bytes, int mode,
                       bemx_server** out) {
    if (!out || nslots < 1 || nslots > 16) return BEMX_ERR_ARG;
    if (mode != BEMX_MODE_FP64 && mode != BEMX_MODE_MIXED) return BEMX_ERR_ARG;
    *out = nullptr;
    if (hipSetDevice(device) != hipSuccess) return BEMX_ERR_HIP;
    size_t freeb = 0, totb = 0;
    if (hipMemGetInfo(&freeb, &totb) != hipSuccess) return BEMX_ERR_HIP;
    auto* s = new bemx_server();
    s->device = device;
    s->mode = mode;
    s->batch = bemx::batch_size();
    s->budget = hbm_budget_bytes > 0 ? (size_t)hbm_budget_bytes : (size_t)(0.8 * (double)freeb);
    // the contexts (and with them the device code of the library) are made
    // on this thread, before any worker starts
    for (int i = 0; i < nslots; ++i) {
        bemx_ctx* c = nullptr;
        int st = bemx_create(device, &c);
        if (st) {
            for (Slot& sl : s->slots) bemx_destroy(sl.ctx);
            delete s;
            return st;
        }
        // a slot's first solve of a mesh size, batched or not, captures the
        // factorisation graphs of every batch size up to the batch limit
        bemx::ctx_capture_ahead(c, s->batch);
        Slot sl;
        sl.ctx = c;
        s->slots.push_back(sl);
    }
    s->slot_captures.assign(nslots, 0);
    s->slot_replays.assign(nslots, 0);
    s->slot_precaptures.assign(nslots, 0);
    for (int i = 0; i < nslots; ++i) s->workers.emplace_back(worker_main, s, i);
    *out = s;
    return BEMX_OK;
}

int bemx_server_submit(bemx_server* s, const double* cx, const double* cy, const double* cz,
                       const int* cone, int np, int n, int nbe, int npg, double ge, double rnu,
                       double dam, double rho, double fr, int nrhs, const double* zdfi,
                       const int* kode, int kode_cols, long long* id) {
    if (!s || !cx || !cy || !cz || !cone || !zdfi || !kode || !id) return BEMX_ERR_ARG;
    if (np <= 0 || n <= 0 || nbe <= 0 || nbe > n || npg <= 0 || nrhs < 1 || nrhs > 4 ||
        !(fr != 0.0) || (kode_cols != 1 && kode_cols != nrhs))
        return BEMX_ERR_ARG;
    auto r = std::make_unique<Request>();
    r->cx.assign(cx, cx + np);
    r->cy.assign(cy, cy + np);
    r->cz.assign(cz, cz + np);
    r->cone.assign(cone, cone + 4 * (size_t)n);
    r->np = np;
    r->n = n;
    r->nbe = nbe;
    r->npg = npg;
    r->ge = ge;
    r->rnu = rnu;
    r->dam = dam;
    r->rho = rho;
    r->fr = fr;
    r->nrhs = nrhs;
    r->zdfi.assign(zdfi, zdfi + (size_t)6 * nbe * nrhs);
    r->kode.assign(kode, kode + (size_t)3 * nbe * kode_cols);
    r->kode_cols = kode_cols;
    uint64_t h = 1469598103934665603ull;
    int dims[4] = {np, n, nbe, npg};
    double mat[2] = {ge, rnu};
    h = fnv1a(h, dims, sizeof dims);
    h = fnv1a(h, mat, sizeof mat);
    h = fnv1a(h, cx, sizeof(double) * np);
    h = fnv1a(h, cy, sizeof(double) * np);
    h = fnv1a(h, cz, sizeof(double) * np);
    h = fnv1a(h, cone, sizeof(int) * 4 * (size_t)n);
    r->mesh_key = h ? h : 1;
    r->footprint = request_footprint(*r, s->mode);
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return BEMX_ERR_STATE;
    if (r->footprint > s->budget) {
        s->err = "request footprint exceeds the HBM budget";
        return BEMX_ERR_NOMEM;
    }
    r->id = s->next_id++;
    r->t_submit = clk::now();
    *id = r->id;
    s->queue.push_back(r.get());
    s->reqs[r->id] = std::move(r);
    s->cv_work.notify_all();
    return BEMX_OK;
}

int bemx_server_wait(bemx_server* s, long long id, double* x, double* info, char* msg,
                     int msglen) {
    if (!s) return BEMX_ERR_ARG;
    std::unique_ptr<Request> r;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        auto it = s->reqs.find(id);
        if (it == s->reqs.end()) return BEMX_ERR_ARG;
        Request* p = it->second.get();
        s->cv_done.wait(lk, [&] { return p->done || s->stop; });
        if (!p->done) return BEMX_ERR_STATE;
        r = std::move(it->second);
        s->reqs.erase(it);
    }
    if (x && !r->status) memcpy(x, r->x.data(), sizeof(double) * r->x.size());
    if (info) {
        info[0] = ms_between(r->t_submit, r->t_done);
        info[1] = ms_between(r->t_submit, r->t_start);
        info[2] = ms_between(r->t_start, r->t_done);
        info[3] = r->stats[0];
        info[4] = r->stats[1];
        info[5] = r->stats[2];
        info[6] = r->slot;
        info[7] = r->stats[3];
        info[8] = ms_between(s->t0, r->t_start);
        info[9] = ms_between(s->t0, r->t_done);
        info[10] = r->batch;
    }
    if (msg && msglen > 0) {
        strncpy(msg, r->err.c_str(), (size_t)msglen - 1);
        msg[msglen - 1] = 0;
    }
    return r->status;
}

int bemx_server_stats(bemx_server* s, double* out, int nout) {
    if (!s || !out || nout < 1) return BEMX_ERR_ARG;
    double v[10];
    {
        std::lock_guard<std::mutex> lk(s->mu);
        v[8] = (double)s->batches;
        v[0] = (double)s->completed;
        v[1] = (double)s->captures;
        v[2] = (double)s->replays;
        v[3] = (double)s->charged;
        v[4] = (double)s->peak_charged;
        v[5] = (double)s->budget;
        v[6] = (double)s->releases;
        v[7] = (double)s->over_charge;
        v[9] = (double)s->precaptures;
    }
    for (int i = 0; i < nout && i < 10; ++i) out[i] = v[i];
    return BEMX_OK;
}

int bemx_server_destroy(bemx_server* s) {
    if (!s) return BEMX_ERR_ARG;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    s->cv_done.notify_all();
    // each worker releases its context's workspace, streams and graphs
    // before it returns
    for (std::thread& t : s->workers) t.join();
    for (Slot& sl : s->slots) bemx_destroy(sl.ctx);
    delete s;
    return BEMX_OK;
}

}  // extern "C"

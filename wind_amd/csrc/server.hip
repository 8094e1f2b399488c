// Request server: continuous batching of integration requests on one GPU.
//
// A background scheduler thread owns `nslots` execution slots (one stream
// each) and a device arena of a fixed HBM budget.  Requests are admitted in
// arrival order when a slot is free and their device footprint fits in the
// arena; each admitted request is copied in, integrated by a persistent
// kernel with a full-device grid, and copied out, all on its slot's stream.
// Admitted requests run concurrently: a request's workgroups exit as its
// queue drains, and the hardware dispatcher hands the freed CUs to the
// workgroups of requests launched after it.  Completion is
// observed with an event; latency runs from submission to results on the
// host.  Resident blocks keep a cell block's state in the arena across
// requests (only results travel); requests on one block run in order, and
// a large block starts each request's systems longest first by their
// sub-step counts in the request before.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/wind.h"
#include "wind_internal.h"

using namespace wind;
using Clock = std::chrono::steady_clock;

namespace {

// First-fit allocator over [0, size) with 256-byte granularity.
class Arena {
   public:
    void init(size_t size) {
        free_.clear();
        free_[0] = size;
        size_ = size;
    }
    bool alloc(size_t bytes, size_t *off) {
        bytes = (bytes + 255) & ~(size_t)255;
        if (bytes == 0) bytes = 256;
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= bytes) {
                *off = it->first;
                const size_t rest = it->second - bytes;
                const size_t roff = it->first + bytes;
                free_.erase(it);
                if (rest) free_[roff] = rest;
                used_ += bytes;
                return true;
            }
        }
        return false;
    }
    void release(size_t off, size_t bytes) {
        bytes = (bytes + 255) & ~(size_t)255;
        if (bytes == 0) bytes = 256;
        used_ -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    size_t used() const { return used_; }
    size_t size() const { return size_; }

   private:
    std::map<size_t, size_t> free_;
    size_t size_ = 0, used_ = 0;
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    unsigned long long *counters = nullptr;  // device, 4 words
    unsigned long long *host_counters = nullptr;  // pinned, 4 words
    char *stage = nullptr;  // pinned staging for a request's host arrays
    size_t stage_bytes = 0;
    size_t out_off = 0, out_bytes = 0;  // where the results sit in `stage`
    int64_t request = -1;
};

}  // namespace

// A block of at least kBlockOrderMin systems also keeps its systems'
// sub-step counts from the last request and a start order: the next
// request on it starts them longest first (results do not depend on the
// order; smaller blocks start every system at once anyway).
constexpr int kBlockOrderMin = 1 << 17;

struct wind_block {
    const wind_network *net = nullptr;
    int nsys = 0, const_stride = 0;
    size_t off = 0, bytes = 0;
    float *constants = nullptr, *equations = nullptr;  // device (in the arena)
    int32_t *steps = nullptr, *order = nullptr;        // device, [nsys] each, or null
    unsigned *count = nullptr;                         // device, kHintCountWords
    bool have_steps = false;
    bool busy = false;
};

namespace {

enum ReqState { PENDING, RUNNING, DONE, FAILED };

struct Request {
    int64_t id = 0;
    wind_request_desc d{};
    Clock::time_point submitted, finished;
    ReqState state = PENDING;
    int slot = -1;
    size_t off = 0, bytes = 0;
    int64_t steps = 0;
    std::string error;
};

}  // namespace

struct wind_server {
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::thread worker;
    bool stop = false;
    char *arena_base = nullptr;
    Arena arena;
    std::vector<Slot> slots;
    std::deque<int64_t> pending;
    std::unordered_map<int64_t, Request> reqs;
    int64_t next_id = 1;
    int num_cu = 0;
    int device = 0;  // the device current when the server was created
    std::string last_error;
    std::vector<wind_block *> blocks;  // live resident blocks (freed by destroy)
};

namespace {

size_t footprint(const Request &r) {
    if (r.d.block) return 0;
    const size_t N = (size_t)r.d.net->N;
    const size_t stride = r.d.const_stride > 0 ? (size_t)r.d.const_stride : (size_t)r.d.net->nconst;
    const size_t cb = ((size_t)(r.d.nsys - 1) * stride + r.d.net->nconst) * sizeof(float);
    return ((cb + 255) & ~(size_t)255) + (size_t)r.d.nsys * N * sizeof(float);
}

// Grows a slot's pinned staging buffer.  Host arrays go through it so that
// copies are truly asynchronous and the scheduler never blocks on a slot.
bool ensure_stage(Slot &s, size_t bytes) {
    if (bytes <= s.stage_bytes) return true;
    if (s.stage) {
        (void)hipStreamSynchronize(s.stream);
        (void)hipHostFree(s.stage);
    }
    s.stage = nullptr;
    s.stage_bytes = 0;
    const size_t want = std::max(bytes, (size_t)1 << 20);
    if (hipHostMalloc((void **)&s.stage, want, 0) != hipSuccess) {
        s.stage = nullptr;
        return false;
    }
    s.stage_bytes = want;
    return true;
}

// Enqueues an admitted request on its slot.  Called with srv->mu held.
bool start(wind_server *srv, Request &r, int si) {
    Slot &s = srv->slots[si];
    const wind_network *net = r.d.block ? r.d.block->net : r.d.net;
    const int nsys = r.d.block ? r.d.block->nsys : r.d.nsys;
    const int stride = r.d.block ? r.d.block->const_stride : (r.d.const_stride > 0 ? r.d.const_stride : net->nconst);
    const size_t cb = ((size_t)(nsys - 1) * stride + net->nconst) * sizeof(float);
    const size_t cba = (cb + 255) & ~(size_t)255;
    const size_t eb = (size_t)nsys * net->N * sizeof(float);
    if ((r.d.equations || !r.d.block) && !ensure_stage(s, cba + eb)) {
        r.error = "cannot allocate pinned staging memory";
        return false;
    }
    float *dc, *de;
    if (r.d.block) {
        dc = r.d.block->constants;
        de = r.d.block->equations;
        r.d.block->busy = true;
    } else {
        dc = (float *)(srv->arena_base + r.off);
        de = (float *)(srv->arena_base + r.off + cba);
        std::memcpy(s.stage, r.d.constants, cb);
        std::memcpy(s.stage + cba, r.d.equations, eb);
        if (hipMemcpyAsync(dc, s.stage, cba + eb, hipMemcpyHostToDevice, s.stream) != hipSuccess) {
            r.error = "copy to device failed";
            return false;
        }
    }
    wind_zero_counters(s.counters, 4, s.stream);
    IntegrateArgs a;
    a.tnow = r.d.tnow;
    a.tend = r.d.tend;
    a.h0 = (r.d.tend - r.d.tnow) / (float)r.d.n_integration_steps;
    a.ABS = r.d.ABSOLUTE;
    a.REL = r.d.RELATIVE;
    a.controller = r.d.controller;
    a.constants = dc;
    a.equations = de;
    a.nsys = nsys;
    a.const_stride = stride;
    a.max_attempts = kDefaultMaxAttempts;
    a.steps_out = nullptr;
    a.status_out = nullptr;
    a.index = nullptr;
    a.counters = s.counters;
    const bool ordered = r.d.block && r.d.block->steps && wind_queue_kernel(net, r.d.mode, WIND_KERNEL_AUTO);
    if (ordered) {
        if (r.d.block->have_steps) {
            wind_order_by_hint(r.d.block->steps, nsys, r.d.block->order, r.d.block->count, s.stream);
            a.index = r.d.block->order;
        }
        a.steps_out = r.d.block->steps;
    }
    // Full-device grid (measured against a fixed 1/nslots share of the
    // device: more requests per second and lower p50 and p99 latency).
    SolverChoice sc;
    if (nsys > 0 && r.d.tnow < r.d.tend) {
        int rc = wind_launch_async(net, a, r.d.method, r.d.mode, WIND_KERNEL_AUTO, sc, s.stream);
        if (rc) {
            r.error = wind_last_error();
            if (r.d.block) r.d.block->busy = false;
            return false;
        }
        if (ordered) r.d.block->have_steps = true;
    }
    s.out_off = cba;
    s.out_bytes = r.d.equations ? eb : 0;
    if (s.out_bytes) (void)hipMemcpyAsync(s.stage + cba, de, eb, hipMemcpyDeviceToHost, s.stream);
    (void)hipMemcpyAsync(s.host_counters, s.counters, 4 * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                         s.stream);
    (void)hipEventRecord(s.done, s.stream);
    s.request = r.id;
    r.slot = si;
    r.state = RUNNING;
    return true;
}

void scheduler(wind_server *srv) {
    (void)hipSetDevice(srv->device);
    std::unique_lock<std::mutex> lk(srv->mu);
    while (!srv->stop || !srv->pending.empty() || std::any_of(srv->slots.begin(), srv->slots.end(),
                                                               [](const Slot &s) { return s.request >= 0; })) {
        bool progress = false;
        // completions
        for (size_t si = 0; si < srv->slots.size(); si++) {
            Slot &s = srv->slots[si];
            if (s.request < 0) continue;
            if (hipEventQuery(s.done) != hipSuccess) continue;
            Request &r = srv->reqs[s.request];
            if (s.out_bytes) std::memcpy(r.d.equations, s.stage + s.out_off, s.out_bytes);
            r.finished = Clock::now();
            r.steps = (int64_t)s.host_counters[1];
            r.state = DONE;
            if (r.d.block)
                r.d.block->busy = false;
            else
                srv->arena.release(r.off, r.bytes);
            s.request = -1;
            progress = true;
        }
        if (progress) srv->cv_done.notify_all();
        // admissions, in arrival order; a request whose block is busy lets later ones pass
        for (auto it = srv->pending.begin(); it != srv->pending.end();) {
            int free_slot = -1;
            for (size_t si = 0; si < srv->slots.size(); si++)
                if (srv->slots[si].request < 0) {
                    free_slot = (int)si;
                    break;
                }
            if (free_slot < 0) break;
            Request &r = srv->reqs[*it];
            if (r.d.block && r.d.block->busy) {
                ++it;
                continue;
            }
            r.bytes = footprint(r);
            if (r.bytes && !srv->arena.alloc(r.bytes, &r.off)) break;  // wait for memory
            if (!start(srv, r, free_slot)) {
                if (r.bytes) srv->arena.release(r.off, r.bytes);
                r.state = FAILED;
                r.finished = Clock::now();
                srv->cv_done.notify_all();
            }
            it = srv->pending.erase(it);
            progress = true;
        }
        if (!progress) srv->cv_work.wait_for(lk, std::chrono::microseconds(50));
    }
}

int sfail(wind_server *srv, int code, const std::string &m) {
    if (srv) srv->last_error = m;
    return wind_fail(code, m.c_str());
}

}  // namespace

extern "C" {

int wind_server_create(size_t hbm_budget_bytes, int nslots, wind_server **out) {
    if (!out) return wind_fail(WIND_ERR_ARGUMENT, "out is NULL");
    *out = nullptr;
    if (nslots < 1 || nslots > 16) return wind_fail(WIND_ERR_ARGUMENT, "nslots must be 1-16");
    if (hbm_budget_bytes < (1u << 20)) return wind_fail(WIND_ERR_ARGUMENT, "HBM budget below 1 MiB");
    int ncu = 0;
    int rc = wind_device_cus(&ncu);
    if (rc) return rc;
    wind_server *srv = new wind_server();
    srv->num_cu = ncu;
    (void)hipGetDevice(&srv->device);
    if (hipMalloc((void **)&srv->arena_base, hbm_budget_bytes) != hipSuccess) {
        delete srv;
        return wind_fail(WIND_ERR_DEVICE, "cannot allocate the server arena");
    }
    srv->arena.init(hbm_budget_bytes);
    srv->slots.resize(nslots);
    for (Slot &s : srv->slots) {
        if (hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&s.done, hipEventDisableTiming) != hipSuccess ||
            hipMalloc((void **)&s.counters, 4 * sizeof(unsigned long long)) != hipSuccess ||
            hipHostMalloc((void **)&s.host_counters, 4 * sizeof(unsigned long long), 0) != hipSuccess) {
            wind_server_destroy(srv);
            return wind_fail(WIND_ERR_DEVICE, "cannot create server slots");
        }
    }
    srv->worker = std::thread(scheduler, srv);
    *out = srv;
    return WIND_OK;
}

void wind_server_destroy(wind_server *srv) {
    if (!srv) return;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->stop = true;
    }
    srv->cv_work.notify_all();
    if (srv->worker.joinable()) srv->worker.join();
    for (Slot &s : srv->slots) {
        if (s.stream) (void)hipStreamSynchronize(s.stream);
        if (s.stream) (void)hipStreamDestroy(s.stream);
        if (s.done) (void)hipEventDestroy(s.done);
        if (s.counters) (void)hipFree(s.counters);
        if (s.host_counters) (void)hipHostFree(s.host_counters);
        if (s.stage) (void)hipHostFree(s.stage);
    }
    for (wind_block *b : srv->blocks) delete b;
    if (srv->arena_base) (void)hipFree(srv->arena_base);
    delete srv;
}

int wind_server_block_create(wind_server *srv, const wind_network *net, const float *constants, int const_stride,
                             const float *equations, int nsys, wind_block **out) {
    if (!srv || !net || !out || nsys < 1 || !constants || !equations)
        return sfail(srv, WIND_ERR_ARGUMENT, "bad block arguments");
    *out = nullptr;
    wind_block *b = new wind_block();
    b->net = net;
    b->nsys = nsys;
    b->const_stride = const_stride > 0 ? const_stride : net->nconst;
    const size_t cb = ((size_t)(nsys - 1) * b->const_stride + net->nconst) * sizeof(float);
    const size_t eb = (size_t)nsys * net->N * sizeof(float);
    const size_t ebp = (eb + 255) & ~(size_t)255;
    const size_t ib = nsys >= kBlockOrderMin ? (((size_t)nsys * 4 + 255) & ~(size_t)255) : 0;
    b->bytes = ((cb + 255) & ~(size_t)255) + ebp + 2 * ib + (ib ? kHintCountWords * sizeof(unsigned) : 0);
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        if (!srv->arena.alloc(b->bytes, &b->off)) {
            delete b;
            return sfail(srv, WIND_ERR_ARGUMENT, "block does not fit in the server's HBM budget");
        }
    }
    b->constants = (float *)(srv->arena_base + b->off);
    b->equations = (float *)(srv->arena_base + b->off + ((cb + 255) & ~(size_t)255));
    if (ib) {
        char *extra = (char *)b->equations + ebp;
        b->steps = (int32_t *)extra;
        b->order = (int32_t *)(extra + ib);
        b->count = (unsigned *)(extra + 2 * ib);
    }
    if (hipMemcpy(b->constants, constants, cb, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->equations, equations, eb, hipMemcpyHostToDevice) != hipSuccess) {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->arena.release(b->off, b->bytes);
        delete b;
        return sfail(srv, WIND_ERR_DEVICE, "block upload failed");
    }
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->blocks.push_back(b);
    }
    *out = b;
    return WIND_OK;
}

int wind_server_block_read(wind_server *srv, wind_block *b, float *equations) {
    if (!srv || !b || !equations) return sfail(srv, WIND_ERR_ARGUMENT, "bad arguments");
    std::unique_lock<std::mutex> lk(srv->mu);
    srv->cv_done.wait(lk, [&] {
        if (b->busy) return false;
        for (int64_t id : srv->pending)
            if (srv->reqs[id].d.block == b) return false;
        return true;
    });
    lk.unlock();
    if (hipMemcpy(equations, b->equations, (size_t)b->nsys * b->net->N * sizeof(float), hipMemcpyDeviceToHost) !=
        hipSuccess)
        return sfail(srv, WIND_ERR_DEVICE, "block read failed");
    return WIND_OK;
}

void wind_server_block_destroy(wind_server *srv, wind_block *b) {
    if (!srv || !b) return;
    std::unique_lock<std::mutex> lk(srv->mu);
    // wait until no running or queued request refers to the block
    srv->cv_done.wait(lk, [&] {
        if (b->busy) return false;
        for (int64_t id : srv->pending)
            if (srv->reqs[id].d.block == b) return false;
        return true;
    });
    srv->arena.release(b->off, b->bytes);
    srv->blocks.erase(std::remove(srv->blocks.begin(), srv->blocks.end(), b), srv->blocks.end());
    delete b;
}

int wind_server_submit(wind_server *srv, const wind_request_desc *d, int64_t *id) {
    if (!srv || !d || !id) return sfail(srv, WIND_ERR_ARGUMENT, "bad arguments");
    if (d->method != WIND_SIE && d->method != WIND_RK2) return sfail(srv, WIND_ERR_ARGUMENT, "bad method");
    if (d->mode != WIND_MODE_EXACT && d->mode != WIND_MODE_FAST) return sfail(srv, WIND_ERR_ARGUMENT, "bad mode");
    if (d->controller != WIND_CTRL_REFERENCE && d->controller != WIND_CTRL_ADAPTIVE)
        return sfail(srv, WIND_ERR_ARGUMENT, "bad controller");
    if (d->n_integration_steps < 0) return sfail(srv, WIND_ERR_ARGUMENT, "n_integration_steps < 0");
    if (!d->block && (!d->net || d->nsys < 1 || !d->constants || !d->equations))
        return sfail(srv, WIND_ERR_ARGUMENT, "a request needs a network and host arrays, or a block");
    std::lock_guard<std::mutex> lk(srv->mu);
    Request r;
    r.id = srv->next_id++;
    r.d = *d;
    r.submitted = Clock::now();
    if (!d->block && footprint(r) > srv->arena.size())
        return sfail(srv, WIND_ERR_ARGUMENT, "request larger than the server's HBM budget");
    srv->reqs[r.id] = r;
    srv->pending.push_back(r.id);
    *id = r.id;
    srv->cv_work.notify_all();
    return WIND_OK;
}

int wind_server_wait(wind_server *srv, int64_t id, double *latency_us, int64_t *total_steps) {
    if (!srv) return wind_fail(WIND_ERR_ARGUMENT, "server is NULL");
    std::unique_lock<std::mutex> lk(srv->mu);
    auto it = srv->reqs.find(id);
    if (it == srv->reqs.end()) return sfail(srv, WIND_ERR_ARGUMENT, "unknown request");
    srv->cv_done.wait(lk, [&] { return it->second.state == DONE || it->second.state == FAILED; });
    Request r = it->second;
    srv->reqs.erase(it);
    lk.unlock();
    if (latency_us) *latency_us = std::chrono::duration<double, std::micro>(r.finished - r.submitted).count();
    if (total_steps) *total_steps = r.steps;
    if (r.state == FAILED) return sfail(srv, WIND_ERR_DEVICE, "request failed: " + r.error);
    return WIND_OK;
}

int wind_server_inflight(wind_server *srv, int *running, int *pending, size_t *arena_used) {
    if (!srv) return wind_fail(WIND_ERR_ARGUMENT, "server is NULL");
    std::lock_guard<std::mutex> lk(srv->mu);
    int run = 0;
    for (const Slot &s : srv->slots) run += s.request >= 0;
    if (running) *running = run;
    if (pending) *pending = (int)srv->pending.size();
    if (arena_used) *arena_used = srv->arena.used();
    return WIND_OK;
}

}  // extern "C"

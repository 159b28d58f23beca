// server.hip — a serving loop over resident meshes (no reference counterpart: the reference answers
// one synchronous call at a time, bvh.cu:80-158).
//
//   * Mesh cache: registered meshes (host arrays, kept by the caller) are uploaded and built on
//     first use and stay resident; when an admission needs room, the least recently used meshes
//     that no in-flight request references are evicted.
//   * Admission by HBM budget: a request is admitted (FIFO) when its device buffer (queries in,
//     distances and triangles out: 20 B per query, plus its own query workspace: the fixup queue
//     and the Morton sort of its points) plus any mesh it needs fit in the budget next to what is
//     resident and in flight; otherwise it waits for completions.
//   * Continuous batching: admitted requests go round-robin onto a small pool of streams (upload,
//     query, download, completion event) as soon as they are admitted, without batch boundaries;
//     a worker thread admits, launches and retires.  Each request works in its own buffer, so
//     requests on the same mesh overlap on the device too.
//   * Meshes load on a stream of their own: a load waits for its build only, not for the requests
//     queued on the serving streams.
//   * Latency = completion (results in host memory) minus arrival (the submit time, or an arrival
//     timestamp on the server clock given by the caller); statistics keep every request's latency.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <list>
#include <mutex>
#include <new>
#include <thread>
#include <unordered_map>
#include <vector>

#include "query.hpp"

namespace ssdf_amd {
// resident-mesh entry points of capi.hip
int server_mesh_create(const int *s0, const int *s1, const int *s2, ptrdiff_t nsel, const float *sx,
                       const float *sy, const float *sz, ptrdiff_t nsp, hipStream_t stream, ssdf_mesh_t *out,
                       size_t *bytes);
// Destroys a mesh of server_mesh_create, its frees ordered on the server's (live) streams.
void server_mesh_destroy(ssdf_mesh_t h);
}  // namespace ssdf_amd

using namespace ssdf_amd;

namespace {

using Clock = std::chrono::steady_clock;

double seconds_since(Clock::time_point t0) {
    return std::chrono::duration<double>(Clock::now() - t0).count();
}

struct MeshEntry {
    const int *s0, *s1, *s2;
    const float *sx, *sy, *sz;
    ptrdiff_t nsel, nsp;
    ssdf_mesh_t handle = nullptr;
    size_t bytes = 0;
    int inflight = 0;
    std::list<int>::iterator lru;  // position in the LRU list while resident
    bool resident = false;
};

struct Request {
    uint64_t ticket;
    int mesh;
    int mode;
    ptrdiff_t n;
    const float *x, *y, *z;
    float *out_d;
    int *out_tri;
    unsigned flags;
    double arrival;
    double done_at = -1.0;
    int rc = 0;
    // in flight
    int slot = -1;
    void *dev = nullptr;  // x, y, z, out_d, out_tri on the device
    size_t bytes = 0;
    hipEvent_t ev = nullptr;
};

constexpr int kStreams = 4;

}  // namespace

struct ssdf_server {
    size_t budget = 0;
    size_t used = 0, peak = 0;
    Clock::time_point t0 = Clock::now();
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::unordered_map<int, MeshEntry> meshes;
    std::list<int> lru;  // front = most recently used
    std::deque<Request *> pending;
    std::vector<Request *> inflight;
    std::unordered_map<uint64_t, Request *> by_ticket;
    uint64_t next_ticket = 1;
    hipStream_t streams[kStreams] = {};
    hipStream_t load_stream = nullptr;  // mesh uploads and builds
    int next_stream = 0;
    int max_inflight = 16;
    bool stop = false;
    std::thread worker;
    int device = 0;
    // statistics
    std::vector<double> latencies;
    unsigned long long served = 0, queries = 0, mesh_loads = 0, evictions = 0, errors = 0;
    double first_arrival = -1.0, last_done = 0.0;
};

namespace {

void touch(ssdf_server *s, int id) {
    MeshEntry &e = s->meshes[id];
    s->lru.erase(e.lru);
    s->lru.push_front(id);
    e.lru = s->lru.begin();
}

// Evicts idle meshes (LRU first) until `need` more bytes fit; true if they do.
bool make_room(ssdf_server *s, size_t need) {
    while (s->used + need > s->budget) {
        int victim = -1;
        for (auto it = s->lru.rbegin(); it != s->lru.rend(); ++it)
            if (s->meshes[*it].inflight == 0) {
                victim = *it;
                break;
            }
        if (victim < 0) return false;
        MeshEntry &e = s->meshes[victim];
        server_mesh_destroy(e.handle);
        e.handle = nullptr;
        e.resident = false;
        s->lru.erase(e.lru);
        s->used -= e.bytes;
        s->evictions++;
    }
    return true;
}

// What a resident mesh is charged: an upper bound of everything it can allocate (arrays, LBVH,
// pseudonormals, build workspace, and the winding-number tree that winding-number signs build on
// first use), so that admission never lets the device memory in use pass the budget.
size_t mesh_charge(const MeshEntry &e) {
    return mesh_bytes_bound(e.nsel, e.nsp, 0, true) + gwn_bytes_bound(e.nsel);
}

// Device bytes of a request: points in, distances and triangles out, then its query workspace.
size_t request_io_bytes(const Request *r) { return round256((size_t)r->n * 20); }
size_t request_bytes_of(const Request *r) {
    return request_io_bytes(r) + query_ws_bytes((size_t)r->n, r->n >= kSortMin);
}

// Admits the head of the queue if it fits; returns false when it has to wait.
bool admit_one(ssdf_server *s, std::unique_lock<std::mutex> &lk) {
    if (s->pending.empty() || (int)s->inflight.size() >= s->max_inflight) return false;
    Request *r = s->pending.front();
    auto mit = s->meshes.find(r->mesh);
    MeshEntry &e = mit->second;
    const size_t rb = request_bytes_of(r);
    const size_t mb = e.resident ? 0 : mesh_charge(e);
    if (!make_room(s, rb + mb)) return false;
    s->pending.pop_front();
    const int slot = s->next_stream;
    s->next_stream = (s->next_stream + 1) % kStreams;
    hipStream_t st = s->streams[slot];
    int rc = SSDF_OK;
    if (!e.resident) {
        // upload and build on the load stream; the build is complete when the call returns
        lk.unlock();
        size_t bytes = 0;
        rc = server_mesh_create(e.s0, e.s1, e.s2, e.nsel, e.sx, e.sy, e.sz, e.nsp, s->load_stream, &e.handle,
                                &bytes);
        lk.lock();
        if (!rc) {
            e.resident = true;
            e.bytes = mb;  // the bound, not what the build has allocated so far
            s->used += e.bytes;
            s->lru.push_front(r->mesh);
            e.lru = s->lru.begin();
            s->mesh_loads++;
        }
    } else {
        touch(s, r->mesh);
    }
    if (!rc) {
        r->bytes = rb;
        s->used += rb;
        s->peak = std::max(s->peak, s->used);
        e.inflight++;
        r->slot = slot;
        const size_t n = (size_t)r->n;
        rc = (int)hipMallocAsync(&r->dev, rb, st) == hipSuccess ? SSDF_OK : SSDF_ERR_OOM;
        char *d = (char *)r->dev;
        float *dx = (float *)d, *dy = dx + n, *dz = dy + n, *dd = dz + n;
        int *dt = (int *)(dd + n);
        if (!rc) {
            lk.unlock();
            hipError_t he = hipMemcpyAsync(dx, r->x, 4 * n, hipMemcpyHostToDevice, st);
            if (he == hipSuccess) he = hipMemcpyAsync(dy, r->y, 4 * n, hipMemcpyHostToDevice, st);
            if (he == hipSuccess) he = hipMemcpyAsync(dz, r->z, 4 * n, hipMemcpyHostToDevice, st);
            // EDF reads the current best and keeps it (and its triangle) where it does not improve
            if (he == hipSuccess && r->mode == SSDF_MODE_EDF)
                he = hipMemcpyAsync(dd, r->out_d, 4 * n, hipMemcpyHostToDevice, st);
            if (he == hipSuccess && r->mode == SSDF_MODE_EDF && r->out_tri)
                he = hipMemcpyAsync(dt, r->out_tri, 4 * n, hipMemcpyHostToDevice, st);
            if (he != hipSuccess) rc = SSDF_ERR_HIP;
            if (!rc)
                rc = ws_query_points(e.handle, d + request_io_bytes(r), false, r->mode, r->n, dx, dy, dz, dd,
                                     r->out_tri ? dt : nullptr, nullptr, nullptr, nullptr, r->flags, st);
            if (!rc) {
                he = hipMemcpyAsync(r->out_d, dd, 4 * n, hipMemcpyDeviceToHost, st);
                if (he == hipSuccess && r->out_tri) he = hipMemcpyAsync(r->out_tri, dt, 4 * n, hipMemcpyDeviceToHost, st);
                if (he == hipSuccess) he = hipFreeAsync(r->dev, st);
                if (he == hipSuccess) r->dev = nullptr;
                if (he == hipSuccess) he = hipEventCreateWithFlags(&r->ev, hipEventDisableTiming);
                if (he == hipSuccess) he = hipEventRecord(r->ev, st);
                if (he != hipSuccess) rc = SSDF_ERR_HIP;
            }
            lk.lock();
        }
    }
    if (rc) {
        // failed before or at launch: finish it now with the error
        if (r->dev) (void)hipFreeAsync(r->dev, st), r->dev = nullptr;
        if (r->bytes) s->used -= r->bytes, r->bytes = 0;
        if (r->slot >= 0) e.inflight--;
        r->rc = rc;
        r->done_at = seconds_since(s->t0);
        s->errors++;
        s->cv_done.notify_all();
        return true;
    }
    s->inflight.push_back(r);
    return true;
}

// Retires completed requests; true if any completed.
bool retire(ssdf_server *s) {
    bool any = false;
    for (size_t i = 0; i < s->inflight.size();) {
        Request *r = s->inflight[i];
        const hipError_t q = hipEventQuery(r->ev);
        if (q == hipErrorNotReady) {
            i++;
            continue;
        }
        r->done_at = seconds_since(s->t0);
        if (q != hipSuccess) {
            r->rc = SSDF_ERR_HIP;
            s->errors++;
        } else {
            s->latencies.push_back(r->done_at - r->arrival);
            s->served++;
            s->queries += (unsigned long long)r->n;
            s->last_done = std::max(s->last_done, r->done_at);
        }
        (void)hipEventDestroy(r->ev);
        r->ev = nullptr;
        s->used -= r->bytes;
        r->bytes = 0;
        s->meshes[r->mesh].inflight--;
        s->inflight[i] = s->inflight.back();
        s->inflight.pop_back();
        any = true;
    }
    if (any) s->cv_done.notify_all();
    return any;
}

void worker_loop(ssdf_server *s) {
    (void)hipSetDevice(s->device);
    std::unique_lock<std::mutex> lk(s->mu);
    while (true) {
        bool progressed = retire(s);
        while (admit_one(s, lk)) progressed = true;
        if (s->stop && s->pending.empty() && s->inflight.empty()) break;
        if (!progressed) {
            if (s->inflight.empty() && s->pending.empty())
                s->cv_work.wait(lk);
            else
                s->cv_work.wait_for(lk, std::chrono::microseconds(50));
        }
    }
}

}  // namespace

extern "C" {

int ssdf_server_create(size_t hbm_budget_bytes, int max_inflight, ssdf_server_t *out) {
    if (!out || hbm_budget_bytes < (1u << 20) || max_inflight < 0 || max_inflight > 1024) {
        set_error("ssdf_server_create: budget >= 1 MiB, 0 <= max_inflight <= 1024");
        return SSDF_ERR_INVALID;
    }
    ssdf_server *s = new (std::nothrow) ssdf_server();
    if (!s) return SSDF_ERR_OOM;
    s->budget = hbm_budget_bytes;
    if (max_inflight > 0) s->max_inflight = max_inflight;
    (void)hipGetDevice(&s->device);
    for (int i = 0; i <= kStreams; i++) {
        hipStream_t *sp = i < kStreams ? &s->streams[i] : &s->load_stream;
        if (hipStreamCreateWithFlags(sp, hipStreamNonBlocking) != hipSuccess) {
            for (int j = 0; j < i; j++) (void)hipStreamDestroy(s->streams[j]);
            delete s;
            set_error("ssdf_server_create: stream creation failed");
            return SSDF_ERR_HIP;
        }
    }
    s->worker = std::thread(worker_loop, s);
    *out = s;
    return SSDF_OK;
}

int ssdf_server_add_mesh(ssdf_server_t s, int mesh_id, ptrdiff_t nselements, const int *s0, const int *s1,
                         const int *s2, ptrdiff_t nspoints, const float *sx, const float *sy, const float *sz) {
    if (!s || nselements <= 0 || nspoints <= 0 || !s0 || !s1 || !s2 || !sx || !sy || !sz) {
        set_error("ssdf_server_add_mesh: invalid arguments");
        return SSDF_ERR_INVALID;
    }
    std::lock_guard<std::mutex> g(s->mu);
    if (s->meshes.count(mesh_id)) {
        set_error("ssdf_server_add_mesh: mesh id already registered");
        return SSDF_ERR_INVALID;
    }
    MeshEntry e{s0, s1, s2, sx, sy, sz, nselements, nspoints};
    s->meshes.emplace(mesh_id, e);
    return SSDF_OK;
}

int ssdf_server_charge(ssdf_server_t s, int mesh_id, ptrdiff_t npoints, size_t *mesh_bytes, size_t *request_bytes) {
    if (!s || npoints < 0) {
        set_error("ssdf_server_charge: invalid arguments");
        return SSDF_ERR_INVALID;
    }
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->meshes.find(mesh_id);
    if (it == s->meshes.end()) {
        set_error("ssdf_server_charge: unknown mesh id");
        return SSDF_ERR_INVALID;
    }
    Request r{};
    r.n = npoints;
    if (mesh_bytes) *mesh_bytes = mesh_charge(it->second);
    if (request_bytes) *request_bytes = request_bytes_of(&r);
    return SSDF_OK;
}

int ssdf_server_clock(ssdf_server_t s, double *now) {
    if (!s || !now) {
        set_error("ssdf_server_clock: null argument");
        return SSDF_ERR_INVALID;
    }
    *now = seconds_since(s->t0);
    return SSDF_OK;
}

int ssdf_server_submit(ssdf_server_t s, int mesh_id, int mode, ptrdiff_t npoints, const float *x, const float *y,
                       const float *z, float *out_d, int *out_tri, unsigned flags, double arrival,
                       uint64_t *ticket) {
    if (!s || !ticket || npoints <= 0 || !x || !y || !z || !out_d ||
        (mode != SSDF_MODE_EDF && mode != SSDF_MODE_SDF)) {
        set_error("ssdf_server_submit: invalid arguments (modes EDF or SDF, at least one point)");
        return SSDF_ERR_INVALID;
    }
    Request *r = new (std::nothrow) Request();
    if (!r) return SSDF_ERR_OOM;
    *r = Request{0, mesh_id, mode, npoints, x, y, z, out_d, out_tri, flags, 0.0};
    {
        std::lock_guard<std::mutex> g(s->mu);
        auto it = s->meshes.find(mesh_id);
        if (it == s->meshes.end()) {
            delete r;
            set_error("ssdf_server_submit: unknown mesh id");
            return SSDF_ERR_INVALID;
        }
        if (request_bytes_of(r) + (it->second.resident ? 0 : mesh_charge(it->second)) > s->budget) {
            delete r;
            set_error("ssdf_server_submit: request larger than the HBM budget");
            return SSDF_ERR_OOM;
        }
        r->arrival = arrival > 0.0 ? arrival : seconds_since(s->t0);
        if (s->first_arrival < 0.0 || r->arrival < s->first_arrival) s->first_arrival = r->arrival;
        r->ticket = s->next_ticket++;
        s->by_ticket[r->ticket] = r;
        s->pending.push_back(r);
        *ticket = r->ticket;
    }
    s->cv_work.notify_one();
    return SSDF_OK;
}

int ssdf_server_wait(ssdf_server_t s, uint64_t ticket, double *latency_s) {
    if (!s) {
        set_error("ssdf_server_wait: null server");
        return SSDF_ERR_INVALID;
    }
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->by_ticket.find(ticket);
    if (it == s->by_ticket.end()) {
        set_error("ssdf_server_wait: unknown ticket");
        return SSDF_ERR_INVALID;
    }
    Request *r = it->second;
    s->cv_done.wait(lk, [r] { return r->done_at >= 0.0; });
    const int rc = r->rc;
    if (latency_s) *latency_s = r->done_at - r->arrival;
    s->by_ticket.erase(it);
    delete r;
    if (rc) set_error("ssdf_server: request failed");
    return rc;
}

int ssdf_server_poll(ssdf_server_t s, uint64_t ticket, int *done) {
    if (!s || !done) {
        set_error("ssdf_server_poll: null argument");
        return SSDF_ERR_INVALID;
    }
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->by_ticket.find(ticket);
    if (it == s->by_ticket.end()) {
        set_error("ssdf_server_poll: unknown ticket");
        return SSDF_ERR_INVALID;
    }
    *done = it->second->done_at >= 0.0 ? 1 : 0;
    return SSDF_OK;
}

int ssdf_server_stats(ssdf_server_t s, ssdf_server_stats_t *st) {
    if (!s || !st) {
        set_error("ssdf_server_stats: null argument");
        return SSDF_ERR_INVALID;
    }
    std::lock_guard<std::mutex> g(s->mu);
    std::memset(st, 0, sizeof(*st));
    st->served = s->served;
    st->queries = s->queries;
    st->errors = s->errors;
    st->mesh_loads = s->mesh_loads;
    st->evictions = s->evictions;
    st->pending = s->pending.size();
    st->inflight = s->inflight.size();
    st->bytes_in_use = s->used;
    st->peak_bytes = s->peak;
    if (!s->latencies.empty()) {
        std::vector<double> l = s->latencies;
        auto pct = [&l](double p) {
            const size_t k = std::min(l.size() - 1, (size_t)(p * (double)(l.size() - 1) + 0.5));
            std::nth_element(l.begin(), l.begin() + (ptrdiff_t)k, l.end());
            return l[k];
        };
        st->p50_s = pct(0.50);
        st->p99_s = pct(0.99);
        st->max_s = *std::max_element(l.begin(), l.end());
        if (s->last_done > s->first_arrival) st->throughput_qps = (double)s->queries / (s->last_done - s->first_arrival);
    }
    return SSDF_OK;
}

int ssdf_server_destroy(ssdf_server_t s) {
    if (!s) return SSDF_OK;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->worker.joinable()) s->worker.join();
    for (auto &kv : s->meshes)
        if (kv.second.handle) server_mesh_destroy(kv.second.handle);
    for (auto &kv : s->by_ticket) delete kv.second;
    for (int i = 0; i < kStreams; i++) (void)hipStreamDestroy(s->streams[i]);
    (void)hipStreamDestroy(s->load_stream);
    delete s;
    return SSDF_OK;
}

int ssdf_host_alloc(void **ptr, size_t bytes) {
    if (!ptr) {
        set_error("ssdf_host_alloc: null pointer");
        return SSDF_ERR_INVALID;
    }
    SSDF_HIP_TRY(hipHostMalloc(ptr, std::max<size_t>(bytes, 1), hipHostMallocDefault));
    return SSDF_OK;
}

int ssdf_host_free(void *ptr) {
    if (ptr) SSDF_HIP_TRY(hipHostFree(ptr));
    return SSDF_OK;
}

}  // extern "C"

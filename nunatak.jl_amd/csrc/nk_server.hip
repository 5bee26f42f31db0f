// nk_server.hip — serving form: continuous batching of velocity solves of uneven grid sizes on one GPU.
//
// Host-side scheduler over the solver's C ABI. Worker threads (the slots) each run one request at a time on the
// HIP stream of the handle serving it, so several solves are on the device at once and the small ones fill the
// CUs that a single small grid leaves idle. Handles are pooled by size class (grid + options): a request reuses an
// idle handle of its class, whose device buffers and captured PCG chunk graph stay resident, or gets a new one
// when the HBM budget allows, after idle handles of other classes are evicted least recently used first.
// Admission is in arrival order: the head of the queue waits until a slot is free and its handle fits.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <list>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/nunatak_hip.h"
#include "nk_internal.hpp"

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
    int64_t id = 0;
    nk_grid grid{};
    nk_physics phys{};
    nk_solver_options opts{};
    std::string key;
    int64_t need = 0;                       // device bytes of a new handle of this class
    std::vector<int32_t> ids;
    std::vector<std::vector<double>> data;
    bool done = false;
    nk_status status = NK_OK;
    std::string err;
    std::vector<double> u, v;
    nk_request_info info{};
};

struct Handle {
    nk_solver *h = nullptr;
    std::string key;
    int64_t bytes = 0;
    uint64_t last_used = 0;
    bool busy = false;
};

std::string class_key(const nk_grid &g, const nk_solver_options &o) {
    std::string k(sizeof(g) + sizeof(o), '\0');
    std::memcpy(&k[0], &g, sizeof(g));
    std::memcpy(&k[sizeof(g)], &o, sizeof(o));
    return k;
}

size_t packed_size(const nk_grid &g, const nk_physics &p, int id) {
    int32_t sh[3];
    nk_internal_field_extents(&g, &p, id, sh);
    return (size_t)sh[0] * (size_t)sh[1] * (size_t)sh[2];
}

}  // namespace

struct nk_server {
    nk_server_config cfg{};
    int64_t budget = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<Request>> queue;
    std::unordered_map<int64_t, std::shared_ptr<Request>> live;   // submitted and not yet collected by wait
    std::list<Handle *> handles;
    int64_t held = 0;                     // bytes of every handle (busy or idle) plus reservations being built
    int64_t next_id = 1;
    uint64_t tick = 0;
    bool stop = false;
    std::vector<std::thread> workers;
    Clock::time_point t0;
    std::string err;

    double now_ms() const { return std::chrono::duration<double, std::milli>(Clock::now() - t0).count(); }

    Handle *idle_of(const std::string &key) {
        Handle *best = nullptr;
        for (Handle *h : handles)
            if (!h->busy && h->key == key && (!best || h->last_used > best->last_used)) best = h;
        return best;
    }
    int64_t idle_bytes() const {
        int64_t b = 0;
        for (const Handle *h : handles) if (!h->busy) b += h->bytes;
        return b;
    }
    // Destroy idle handles, least recently used first, until `need` more bytes fit. Caller holds mu.
    void evict_for(int64_t need) {
        while (held + need > budget) {
            auto victim = handles.end();
            for (auto it = handles.begin(); it != handles.end(); ++it)
                if (!(*it)->busy && (victim == handles.end() || (*it)->last_used < (*victim)->last_used)) victim = it;
            if (victim == handles.end()) return;
            Handle *h = *victim;
            handles.erase(victim);
            held -= h->bytes;
            nk_destroy(h->h);
            delete h;
        }
    }

    void finish(const std::shared_ptr<Request> &r) {
        r->done = true;
        cv_done.notify_all();
        cv_work.notify_all();
    }

    nk_status run(Request &r, Handle &hd, bool reused) {
        nk_solver *h = hd.h;
        nk_status st = nk_set_physics(h, &r.phys);
        if (st != NK_OK) return st;
        if (reused) {
            uint8_t keep[NK_FIELD_COUNT] = {0};
            for (int32_t id : r.ids) keep[id] = 1;
            if ((st = nk_internal_reset_fields(h, keep)) != NK_OK) return st;
        }
        bool temperature = false;
        for (size_t k = 0; k < r.ids.size(); k++) {
            if ((st = nk_set_field(h, (nk_field_id)r.ids[k], r.data[k].data(), 0, 0)) != NK_OK) return st;
            temperature |= r.ids[k] == NK_FIELD_TEMPERATURE;
        }
        if (temperature && (st = nk_rate_factor_from_temperature(h)) != NK_OK) return st;
        r.info.start_ms = now_ms();
        nk_info info{};
        if ((st = nk_update_velocities(h, 1, &info)) != NK_OK) return st;
        r.info.picard_iterations = info.picard_iterations;
        r.info.total_linear_iterations = (int32_t)info.total_linear_iterations;
        r.u.resize(packed_size(r.grid, r.phys, NK_FIELD_U));
        r.v.resize(packed_size(r.grid, r.phys, NK_FIELD_V));
        if ((st = nk_get_field(h, NK_FIELD_U, r.u.data(), 0, 0)) != NK_OK) return st;
        if ((st = nk_get_field(h, NK_FIELD_V, r.v.data(), 0, 0)) != NK_OK) return st;
        return NK_OK;
    }

    void worker() {
        for (;;) {
            std::shared_ptr<Request> r;
            Handle *hd = nullptr;
            bool reused = false;
            {
                std::unique_lock<std::mutex> lk(mu);
                for (;;) {
                    if (queue.empty()) {
                        if (stop) return;
                        cv_work.wait(lk);
                        continue;
                    }
                    std::shared_ptr<Request> head = queue.front();
                    if ((hd = idle_of(head->key))) { reused = true; break; }
                    if (head->need > budget) {         // can never fit: fail it rather than block the queue
                        queue.pop_front();
                        head->status = NK_OUT_OF_MEMORY;
                        head->err = "request needs more device memory than the server's HBM budget";
                        head->info.admit_ms = head->info.start_ms = head->info.finish_ms = now_ms();
                        finish(head);
                        continue;
                    }
                    if (held - idle_bytes() + head->need <= budget) {
                        evict_for(head->need);
                        break;
                    }
                    cv_work.wait(lk);
                }
                r = queue.front();
                queue.pop_front();
                r->info.admit_ms = now_ms();
                if (hd) {
                    hd->busy = true;
                } else {
                    held += r->need;                     // reservation while the handle is built
                }
                if (!queue.empty()) cv_work.notify_one();
            }
            nk_status st = NK_OK;
            if (!hd) {
                nk_solver *h = nullptr;
                st = nk_create(&r->grid, &r->phys, &r->opts, &h);
                std::lock_guard<std::mutex> lk(mu);
                if (st != NK_OK) {
                    held -= r->need;
                    r->err = h ? nk_last_error(h) : "nk_create failed";
                    if (h) nk_destroy(h);
                } else {
                    hd = new Handle();
                    hd->h = h;
                    hd->key = r->key;
                    int64_t b = r->need;
                    nk_device_bytes(h, &b);
                    hd->bytes = b;
                    held += b - r->need;
                    hd->busy = true;
                    handles.push_back(hd);
                }
            }
            if (st == NK_OK) {
                st = run(*r, *hd, reused);
                if (st != NK_OK) r->err = nk_last_error(hd->h);
            }
            std::lock_guard<std::mutex> lk(mu);
            r->status = st;
            r->info.status = st;
            r->info.reused_handle = reused ? 1 : 0;
            r->info.finish_ms = now_ms();
            if (hd) {
                r->info.device_bytes = hd->bytes;
                hd->busy = false;
                hd->last_used = ++tick;
            }
            finish(r);
        }
    }
};

nk_status nk_server_create(const nk_server_config *config, nk_server **out) {
    if (!config || !out || config->slots < 1 || config->hbm_budget_bytes < 0) return NK_INVALID_ARGUMENT;
    *out = nullptr;
    nk_server *s = new nk_server();
    s->cfg = *config;
    s->budget = config->hbm_budget_bytes;
    if (s->budget == 0) {
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) { delete s; return NK_DEVICE_ERROR; }
        s->budget = (int64_t)(0.8 * (double)free_b);
    }
    s->t0 = Clock::now();
    for (int k = 0; k < config->slots; k++) s->workers.emplace_back([s] { s->worker(); });
    *out = s;
    return NK_OK;
}

nk_status nk_server_destroy(nk_server *s) {
    if (!s) return NK_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (std::thread &t : s->workers) t.join();
    for (Handle *h : s->handles) {
        nk_destroy(h->h);
        delete h;
    }
    delete s;
    return NK_OK;
}

const char *nk_server_last_error(const nk_server *s) { return s ? s->err.c_str() : "null server"; }

nk_status nk_server_submit(nk_server *s, const nk_grid *grid, const nk_physics *physics, const nk_solver_options *options,
                           int32_t nfields, const int32_t *field_ids, const double *const *fields, int64_t *request_id) {
    if (!s || !grid || !physics || !options || !request_id || nfields < 0 || (nfields > 0 && (!field_ids || !fields)))
        return NK_INVALID_ARGUMENT;
    auto r = std::make_shared<Request>();
    r->grid = *grid;
    r->phys = *physics;
    r->opts = *options;
    int64_t need = 0;
    nk_status st = nk_estimate_device_bytes(grid, physics, options, &need);
    if (st != NK_OK) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->err = "nk_server_submit: invalid grid, physics or options";
        return st;
    }
    r->need = need;
    r->key = class_key(*grid, *options);
    for (int32_t k = 0; k < nfields; k++) {
        const int32_t id = field_ids[k];
        if (id < 0 || id >= NK_FIELD_COUNT || !fields[k]) {
            std::lock_guard<std::mutex> lk(s->mu);
            s->err = "nk_server_submit: bad field id or null field";
            return NK_INVALID_ARGUMENT;
        }
        const size_t n = packed_size(*grid, *physics, id);
        r->ids.push_back(id);
        r->data.emplace_back(fields[k], fields[k] + n);
    }
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop) return NK_INVALID_ARGUMENT;
        r->id = s->next_id++;
        r->info.id = r->id;
        r->info.submit_ms = s->now_ms();
        s->live[r->id] = r;
        s->queue.push_back(r);
        *request_id = r->id;
    }
    s->cv_work.notify_one();
    return NK_OK;
}

nk_status nk_server_poll(nk_server *s, int64_t id, int32_t *done) {
    if (!s || !done) return NK_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->live.find(id);
    if (it == s->live.end()) return NK_INVALID_ARGUMENT;
    *done = it->second->done ? 1 : 0;
    return NK_OK;
}

nk_status nk_server_wait(nk_server *s, int64_t id, double *u, double *v, nk_request_info *info) {
    if (!s) return NK_INVALID_ARGUMENT;
    std::shared_ptr<Request> r;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        auto it = s->live.find(id);
        if (it == s->live.end()) { s->err = "nk_server_wait: unknown request id"; return NK_INVALID_ARGUMENT; }
        r = it->second;
        s->cv_done.wait(lk, [&] { return r->done; });
        s->live.erase(id);
        if (r->status != NK_OK) s->err = r->err;
    }
    if (r->status == NK_OK) {
        if (u) std::copy(r->u.begin(), r->u.end(), u);
        if (v) std::copy(r->v.begin(), r->v.end(), v);
    }
    if (info) *info = r->info;
    return r->status;
}

// Continuous-batching request loop over the engine. A request is a whole
// simulation (initial state, parameters, step count). Requests are admitted
// in arrival order while the HBM budget and the active-slot limit allow;
// every tick co-batches the steps of all admitted simulations of one
// (kind, precision) into shared launches, retires finished ones and frees
// their memory for the queue. A request whose simulation fails on the device
// (non-finite state, unsupported coincidence) fails alone.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <vector>

#include "pg_engine.h"
#include "physim_gpu.h"

namespace {

struct Request {
    uint64_t id = 0;
    std::vector<Entity> state;  // input, then result
    pg_params params{};
    uint32_t steps = 0;
    uint32_t remaining = 0;
    pg_sim* sim = nullptr;
    pg_request_info info{};
    bool done = false;
};

}  // namespace

struct pg_server {
    pg_engine* eng = nullptr;
    uint32_t max_active = 0;
    std::mutex mu;
    uint64_t next_id = 1;
    std::deque<Request*> queue;
    std::vector<Request*> active;
    std::map<uint64_t, Request*> all;
    std::chrono::steady_clock::time_point t0;

    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
};

extern "C" int pg_server_create(pg_engine* eng, uint32_t max_active, pg_server** out) {
    if (!eng || !out) return pg_set_error(PG_ERR_ARG, "engine or out is NULL");
    pg_server* s = new pg_server();
    s->eng = eng;
    s->max_active = max_active;
    s->t0 = std::chrono::steady_clock::now();
    *out = s;
    return PG_OK;
}

extern "C" void pg_server_destroy(pg_server* s) {
    if (!s) return;
    for (auto& kv : s->all) {
        if (kv.second->sim) pg_sim_destroy(kv.second->sim);
        delete kv.second;
    }
    delete s;
}

extern "C" double pg_server_clock(pg_server* s) { return s ? s->now() : 0.0; }

extern "C" int pg_server_submit(pg_server* s, const Entity* state, uint64_t n,
                                const pg_params* params, uint32_t steps, uint64_t* request_id) {
    if (!s || !params || (n && !state)) return pg_set_error(PG_ERR_ARG, "NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    Request* r = new Request();
    r->id = s->next_id++;
    r->state.assign(state, state + n);
    r->params = *params;
    r->steps = r->remaining = steps;
    r->info.id = r->id;
    r->info.n = n;
    r->info.steps = steps;
    r->info.status = PG_OK;
    r->info.submit_s = s->now();
    r->info.admit_s = -1.0;
    r->info.complete_s = -1.0;
    s->queue.push_back(r);
    s->all[r->id] = r;
    if (request_id) *request_id = r->id;
    return PG_OK;
}

static void finish(pg_server* s, Request* r, uint32_t status) {
    if (status == PG_OK && r->sim)
        status = (uint32_t)pg_sim_download(r->sim, r->state.data(), r->state.size());
    if (r->sim) pg_sim_destroy(r->sim);
    r->sim = nullptr;
    r->info.status = status;
    r->info.complete_s = s->now();
    r->done = true;
}

extern "C" int pg_server_tick(pg_server* s, uint32_t max_steps, uint32_t* n_active,
                              uint32_t* n_queued) {
    if (!s) return pg_set_error(PG_ERR_ARG, "server is NULL");
    std::lock_guard<std::mutex> lk(s->mu);
    if (max_steps == 0) max_steps = 1;
    // admission, in arrival order
    while (!s->queue.empty() && (s->max_active == 0 || s->active.size() < s->max_active)) {
        Request* r = s->queue.front();
        uint64_t used = 0, budget = 0;
        pg_engine_memory(s->eng, &used, &budget);
        const uint64_t need = pg_sim_bytes(r->state.size(), r->params.kind);
        if (!s->active.empty() && used + need > budget) break;  // wait for memory
        s->queue.pop_front();
        r->info.admit_s = s->now();
        int rc = pg_sim_create(s->eng, r->state.size(), &r->params, &r->sim);
        if (rc == PG_OK) rc = pg_sim_upload(r->sim, r->state.data(), r->state.size());
        if (rc != PG_OK) {
            finish(s, r, (uint32_t)rc);
            continue;
        }
        if (r->remaining == 0) {
            finish(s, r, PG_OK);
            continue;
        }
        s->active.push_back(r);
    }
    // co-batched steps per (kind, precision) group
    std::map<int, std::vector<Request*>> groups;
    for (Request* r : s->active) groups[r->params.kind * 4 + r->params.precision].push_back(r);
    for (auto& kv : groups) {
        std::vector<Request*>& g = kv.second;
        uint32_t k = max_steps;
        for (Request* r : g) k = std::min(k, r->remaining);
        std::vector<pg_sim*> sims;
        for (Request* r : g) sims.push_back(r->sim);
        const int rc = pg_batch_step(s->eng, sims.data(), (uint32_t)sims.size(), k);
        for (Request* r : g) {
            pg_stats st;
            const bool failed_alone = pg_sim_stats(r->sim, &st) == PG_OK && st.error != PG_OK;
            if (failed_alone) {
                finish(s, r, st.error);
            } else if (rc != PG_OK && rc != PG_ERR_NONFINITE && rc != PG_ERR_COINCIDENT &&
                       rc != PG_ERR_DEPTH) {
                finish(s, r, (uint32_t)rc);  // a batch-wide failure fails every member
            } else {
                r->remaining -= k;
                if (r->remaining == 0) finish(s, r, PG_OK);
            }
        }
    }
    s->active.erase(std::remove_if(s->active.begin(), s->active.end(),
                                   [](Request* r) { return r->done; }),
                    s->active.end());
    if (n_active) *n_active = (uint32_t)s->active.size();
    if (n_queued) *n_queued = (uint32_t)s->queue.size();
    return PG_OK;
}

extern "C" int pg_server_info(pg_server* s, uint64_t id, pg_request_info* info) {
    if (!s || !info) return pg_set_error(PG_ERR_ARG, "NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(id);
    if (it == s->all.end()) return pg_set_error(PG_ERR_ARG, "unknown request %llu",
                                                (unsigned long long)id);
    *info = it->second->info;
    return it->second->done ? PG_OK : PG_ERR_STATE;
}

extern "C" int pg_server_take(pg_server* s, uint64_t id, Entity* out, uint64_t n,
                              pg_request_info* info) {
    if (!s) return pg_set_error(PG_ERR_ARG, "server is NULL");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(id);
    if (it == s->all.end())
        return pg_set_error(PG_ERR_ARG, "unknown request %llu", (unsigned long long)id);
    Request* r = it->second;
    if (!r->done) return pg_set_error(PG_ERR_STATE, "request %llu not finished",
                                      (unsigned long long)id);
    if (info) *info = r->info;
    if (out) {
        if (n != r->state.size()) return pg_set_error(PG_ERR_ARG, "n mismatch");
        memcpy(out, r->state.data(), n * sizeof(Entity));
    }
    const uint32_t status = r->info.status;
    s->all.erase(it);
    delete r;
    return status == PG_OK ? PG_OK : pg_set_error((int)status, "request failed with status %u", status);
}
